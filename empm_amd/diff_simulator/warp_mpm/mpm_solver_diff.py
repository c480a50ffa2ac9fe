"""MPMWARPDiff with the reference's method names and argument meaning
(phys_sim/diff_simulator/warp_mpm/mpm_solver_diff.py:13-1410), running the
substep on libempm instead of Warp kernels.

Grid post-processes keep the reference's "plugin" shape: every add_* call
appends one mutable parameter object to ``collider_params``, one entry to
``grid_postprocess`` and one host ``modify_bc`` callback (or None), in
registration order.  They are evaluated on the device from a table that is
rebuilt from ``collider_params`` at every substep, so callers may mutate
.point / .velocity / .active between substeps exactly as before.
"""
from __future__ import annotations

import math
import time as _time

import numpy as np
import torch

from ... import runtime as rt
from ..._lib import Collider
from .mpm_data_structure import Dirichlet_collider, GridCollider, MPMModelStruct, MPMStateStruct, friction_alpha

MATERIALS = {"jelly": 0, "metal": 1, "sand": 2, "foam": 3, "snow": 4, "plasticine": 5, "neo-hookean": 6,
             "neo-hookean-von-mises": 7}

KIND_SURFACE, KIND_CUBOID, KIND_BBOX, KIND_GEOMETRY, KIND_MOVER, KIND_MASK = range(6)
GEO_TYPES = {"cube": 1, "sphere": 2, "cylinder": 3, "halfspace": 4, "plane": 4}


def _v3(v):
    if np.ndim(v) == 0:
        return (float(v),) * 3  # a scalar assigned to a vec3 field broadcasts
    v = list(v)
    return (float(v[0]), float(v[1]), float(v[2]))


def _surface_type(surface):
    return {"sticky": 0, "slip": 1, "cut": 11}.get(surface, 2)


def _normalized(v):
    s = 1.0 / math.sqrt(float(sum(float(x) ** 2 for x in v)))
    return [s * float(x) for x in v]


def _to_struct(kind, geo_type, p) -> Collider:
    c = Collider()
    c.kind = kind
    c.geo_type = geo_type
    if kind == KIND_MASK:
        c.mask = rt._ptr(p.mask)
        return c
    c.surface_type = int(p.surface_type)
    c.reset = int(p.reset)
    c.active = 1 if p.active else 0
    c.start_time = float(p.start_time)
    c.end_time = float(p.end_time)
    for f in ("point", "normal", "size", "velocity", "rot_axis"):
        v = _v3(getattr(p, f))
        arr = getattr(c, f)
        arr[0], arr[1], arr[2] = v
    c.angular_velocity = float(p.angular_velocity)
    c.radius = float(p.radius)
    c.half_height = float(p.half_height)
    c.friction = float(p.friction)
    return c


def _advance_point(time, dt, param):
    """Host-side motion of cuboids and geometry movers
    (mpm_solver_diff.py:795-801,860-866): float32 point += dt * velocity
    while start_time <= time < end_time."""
    st = float(np.float32(param.start_time))
    en = float(np.float32(param.end_time))
    if time >= st and time < en:
        p = _v3(param.point)
        v = _v3(param.velocity)
        param.point = tuple(float(np.float32(float(np.float32(p[i])) + dt * float(np.float32(v[i])))) for i in range(3))


class MPMWARPDiff:
    """mpm_solver_diff.py:13-1410"""

    def __init__(self):
        self.time_profile = {}
        self.step_callback = None
        # reverse-mode memory policy of the autograd Functions: 1 records every
        # substep (no recompute); K > 1 keeps a state every K substeps and
        # replays each segment in the backward
        self.checkpoint_every = 1
        # gradient convention of the autograd Functions: "reference_compat"
        # reproduces the reference Tape's (grid-mass adjoint dropped, Trail
        # frames seeded at the last state, Clean seeded by the position
        # only); "exact" is the derivative of the discrete rollout
        self.adjoint_mode = "reference_compat"
        self._ctx = None
        self._ctx_key = None

    # mpm_solver_diff.py:18-33
    def initialize(self, n_particles, n_grid=100, grid_lim=1.0, device="cuda:0"):
        self.n_particles = n_particles
        self.n_grid = n_grid
        self.grid_lim = grid_lim
        self.device = device
        self.time = 0.0
        self.grid_postprocess = []
        self.collider_params = []
        self.modify_bc = []
        self.pre_p2g_operations = []
        self.impulse_params = []
        self.particle_velocity_modifiers = []
        self.particle_velocity_modifier_params = []
        self.step_callback = None

    # ---------------------------------------------------------------- device
    def context(self, n_grid, n_particles, device):
        dev = torch.device(device)
        key = (int(n_grid), str(dev))
        if self._ctx is None or self._ctx_key != key or self._ctx.max_particles < n_particles:
            self._ctx = rt.Context(n_grid, max(int(n_particles), 1), dev)
            self._ctx_key = key
        return self._ctx

    def model_arrays(self, mpm_model: MPMModelStruct, mpm_state: MPMStateStruct) -> rt.ModelArrays:
        n = mpm_state.n_particles
        if mpm_model.yield_stress is None or mpm_model.yield_stress.shape[0] != n:
            raise ValueError("model arrays are not initialised for this state (call MPMModelStruct.init)")
        sel = mpm_state.particle_selection
        if sel is not None and not bool(torch.any(sel != 0)):
            sel = None
        g = mpm_model.gravitational_accelaration
        return rt.ModelArrays(material=int(mpm_model.material), n_grid=int(mpm_model.n_grid),
                              grid_lim=float(mpm_model.grid_lim), dx=float(mpm_model.dx),
                              inv_dx=float(mpm_model.inv_dx), mu=mpm_model.mu, lam=mpm_model.lam,
                              yield_stress=mpm_model.yield_stress, mass=mpm_state.particle_mass,
                              vol=mpm_state.particle_vol, selection=sel,
                              rpic_damping=float(mpm_model.rpic_damping),
                              grid_v_damping_scale=float(mpm_model.grid_v_damping_scale),
                              gravity=_v3(g), alpha=float(mpm_model.alpha), hardening=float(mpm_model.hardening),
                              xi=float(mpm_model.xi), plastic_viscosity=float(mpm_model.plastic_viscosity),
                              softening=float(mpm_model.softening))

    def collider_table(self):
        return [_to_struct(kind, geo, p) for (kind, geo), p in zip(self.grid_postprocess, self.collider_params)]

    def sync_colliders_from(self, structs):
        """Write device-advanced collider points back into collider_params."""
        for (kind, _), p, s in zip(self.grid_postprocess, self.collider_params, structs):
            if kind in (KIND_CUBOID, KIND_MOVER):
                p.point = (float(s.point[0]), float(s.point[1]), float(s.point[2]))

    # ---------------------------------------------------------------- params
    # mpm_solver_diff.py:35-36
    def set_parameters(self, device="cuda:0", **kwargs):
        self.set_parameters_dict(self._model, self._state, kwargs, device)

    # mpm_solver_diff.py:38-104
    def set_parameters_dict(self, mpm_model, mpm_state, kwargs={}, device="cuda:0"):
        self._model, self._state = mpm_model, mpm_state
        if "material" in kwargs:
            if kwargs["material"] not in MATERIALS:
                raise TypeError("Undefined material type")
            mpm_model.material = MATERIALS[kwargs["material"]]
        if "yield_stress" in kwargs:
            mpm_model.yield_stress.fill_(float(kwargs["yield_stress"]))
        if "hardening" in kwargs:
            mpm_model.hardening = kwargs["hardening"]
        if "xi" in kwargs:
            mpm_model.xi = kwargs["xi"]
        if "friction_angle" in kwargs:
            mpm_model.friction_angle = kwargs["friction_angle"]
            mpm_model.alpha = friction_alpha(mpm_model.friction_angle)
        if "rpic_damping" in kwargs:
            mpm_model.rpic_damping = kwargs["rpic_damping"]
        if "plastic_viscosity" in kwargs:
            mpm_model.plastic_viscosity = kwargs["plastic_viscosity"]
        if "softening" in kwargs:
            mpm_model.softening = kwargs["softening"]
        if "grid_v_damping_scale" in kwargs:
            mpm_model.grid_v_damping_scale = kwargs["grid_v_damping_scale"]

    # mpm_solver_diff.py:106-135
    def set_E_nu(self, mpm_model, E, nu, device="cuda:0"):
        if isinstance(E, float):
            mpm_model.E.fill_(E)
        else:
            mpm_model.E.copy_(torch.as_tensor(E, dtype=torch.float32).to(mpm_model.E.device).reshape(-1))
        if isinstance(nu, float):
            mpm_model.nu.fill_(nu)
        else:
            mpm_model.nu.copy_(torch.as_tensor(nu, dtype=torch.float32).to(mpm_model.nu.device).reshape(-1))

    # mpm_solver_diff.py:137-154
    def set_E_nu_from_torch(self, mpm_model, E, nu, device="cuda:0"):
        E_inp = E.item() if E.ndim == 0 else E.detach()
        nu_inp = nu.item() if nu.ndim == 0 else nu.detach()
        self.set_E_nu(mpm_model, E_inp, nu_inp, device=device)

    # mpm_solver_diff.py:156-163
    def prepare_mu_lam(self, mpm_model, mpm_state, device="cuda:0"):
        rt.mu_lam(mpm_model.E, mpm_model.nu, mpm_model.mu, mpm_model.lam)

    # ---------------------------------------------------------------- substep
    def _check_modifiers(self):
        if self.pre_p2g_operations or self.particle_velocity_modifiers:
            raise NotImplementedError("particle-side modifiers (impulse / velocity modifiers) are not supported")

    # mpm_solver_diff.py:165-291
    def p2g2p_differentiable(self, mpm_model, mpm_state, next_state, dt, device="cuda:0"):
        self._check_modifiers()
        n = mpm_state.n_particles
        ctx = self.context(mpm_model.n_grid, n, mpm_state.particle_x.device)
        model = self.model_arrays(mpm_model, mpm_state)
        cols = self.collider_table()
        t0 = _time.perf_counter()
        rt.substep_fwd(ctx, model, cols, self.time, dt, mpm_state.particles(with_fe=True),
                       next_state.particles(with_fe=False))
        for k in range(len(self.grid_postprocess)):
            if self.modify_bc[k] is not None:
                self.modify_bc[k](self.time, dt, self.collider_params[k])
        self.time_profile.setdefault("substep_enqueue", []).append(_time.perf_counter() - t0)
        self.time = self.time + dt

    # mpm_solver_diff.py:293-429 (in-place, non-differentiable variant)
    def p2g2p(self, mpm_model, mpm_state, step, dt, device="cuda:0"):
        nxt = mpm_state.partial_clone(device=device, requires_grad=False)
        self.p2g2p_differentiable(mpm_model, mpm_state, nxt, dt, device)
        for f in ("particle_x", "particle_v", "particle_C", "particle_F_trial"):
            getattr(mpm_state, f).copy_(getattr(nxt, f))

    # mpm_solver_diff.py:431-434
    def print_time_profile(self):
        print("MPM Time profile:")
        for key, value in self.time_profile.items():
            print(key, sum(value))

    # ---------------------------------------------------------------- colliders
    def _register(self, kind, geo_type, param, modify):
        self.collider_params.append(param)
        self.grid_postprocess.append((kind, geo_type))
        self.modify_bc.append(modify)

    # mpm_solver_diff.py:437-531
    def add_surface_collider(self, point, normal, surface="sticky", friction=0.0, start_time=0.0, end_time=999.0):
        point = list(point)
        normal = _normalized(normal)
        p = Dirichlet_collider()
        p.start_time = start_time
        p.end_time = end_time
        p.point = _v3(point)
        p.normal = _v3(normal)
        if surface == "sticky" and friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        p.surface_type = _surface_type(surface)
        p.friction = friction
        self._register(KIND_SURFACE, 0, p, None)

    # mpm_solver_diff.py:533-683
    def add_geometry_collider(self, geo_args: dict):
        p = Dirichlet_collider()
        surface = geo_args.get("surface", "sticky")
        p.active = True
        p.start_time = geo_args.get("start_time", 0.0)
        p.end_time = geo_args.get("end_time", 999.0)
        p.friction = geo_args.get("friction", 0.0)
        if "velocity" in geo_args:
            p.velocity = _v3(geo_args["velocity"])
        if surface == "sticky" and p.friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        p.surface_type = _surface_type(surface)
        geo = self._geometry(geo_args, p)
        self._register(KIND_GEOMETRY, geo, p, None)

    def _geometry(self, geo_args, p):
        if "shape" not in geo_args:
            raise ValueError("geo_args must contain 'shape' key.")
        shape = geo_args["shape"]
        if shape == "cube":
            p.point = _v3(geo_args["center"])
            p.size = _v3(geo_args["size"])
        elif shape == "sphere":
            p.point = _v3(geo_args["center"])
            p.radius = float(geo_args["radius"])
        elif shape == "cylinder":
            p.point = _v3(geo_args["center"])
            p.normal = _v3(_normalized(geo_args["axis"]))
            p.radius = float(geo_args["radius"])
            p.half_height = float(geo_args["half_height"])
        elif shape in ("halfspace", "plane"):
            p.point = _v3(geo_args["point"])
            p.normal = _v3(_normalized(geo_args["normal"]))
        else:
            raise ValueError(f"Unknown geometry type: {shape}")
        return GEO_TYPES[shape]

    # mpm_solver_diff.py:685-804
    def set_grid_velocity_on_geometry(self, geo_args: dict):
        p = Dirichlet_collider()
        p.active = True
        p.start_time = geo_args.get("start_time", 0.0)
        p.end_time = geo_args.get("end_time", 999.0)
        p.velocity = _v3(geo_args.get("velocity", (0.0, 0.0, 0.0)))
        p.angular_velocity = geo_args.get("angular_velocity", 0.0)
        p.rot_axis = _v3(geo_args.get("rot_axis", (1.0, 0.0, 0.0)))
        p.reset = geo_args.get("reset", 0)
        geo = self._geometry(geo_args, p)
        self._register(KIND_MOVER, geo, p, _advance_point)

    # mpm_solver_diff.py:814-869 (a scalar size broadcasts to the three axes)
    def set_velocity_on_cuboid(self, point, size, velocity, start_time=0.0, end_time=999.0, reset=0):
        p = Dirichlet_collider()
        p.start_time = start_time
        p.end_time = end_time
        p.point = _v3(list(point))
        p.size = size
        p.velocity = _v3(velocity)
        p.reset = reset
        self._register(KIND_CUBOID, 0, p, _advance_point)

    # mpm_solver_diff.py:871-938
    def add_bounding_box(self, start_time=0.0, end_time=999.0):
        p = Dirichlet_collider()
        p.start_time = start_time
        p.end_time = end_time
        self._register(KIND_BBOX, 0, p, None)

    # mpm_solver_diff.py:1264-1289
    def enforce_grid_velocity_by_mask(self, selection_mask: torch.Tensor):
        p = GridCollider()
        p.mask = selection_mask.to(torch.int32).contiguous()
        self._register(KIND_MASK, 0, p, None)

    # ---------------------------------------------------------------- particle modifiers
    # mpm_solver_diff.py:943-1212,1291-1410: no reference call path uses them
    def add_impulse_on_particles(self, *args, **kwargs):
        raise NotImplementedError("add_impulse_on_particles is not supported by the HIP path")

    def enforce_particle_velocity_translation(self, *args, **kwargs):
        raise NotImplementedError("enforce_particle_velocity_translation is not supported by the HIP path")

    def enforce_particle_velocity_rotation(self, *args, **kwargs):
        raise NotImplementedError("enforce_particle_velocity_rotation is not supported by the HIP path")

    def enforce_particle_velocity_by_mask(self, *args, **kwargs):
        raise NotImplementedError("enforce_particle_velocity_by_mask is not supported by the HIP path")

    def set_particle_mover(self, *args, **kwargs):
        raise NotImplementedError("set_particle_mover is not supported by the HIP path")

    # mpm_solver_diff.py:1214-1262 (reached through static_pos, which the
    # reference's own callers leave at None)
    def restart_and_compute_F_C(self, mpm_model, mpm_state, target_pos, device):
        """F_trial of every selected particle from the displacement to
        target_pos (set_F_C_p2g, normalisation with dt = 0, set_F_C_g2p:
        mpm_utils.py:843-919), then particle_x = target_pos for all particles
        (set_vec3_to_vec3).  C is left as it is, as in the reference."""
        n = mpm_state.n_particles
        x = mpm_state.particle_x
        ctx = self.context(mpm_model.n_grid, n, x.device)
        model = self.model_arrays(mpm_model, mpm_state)
        target = torch.as_tensor(target_pos).to(device=x.device, dtype=torch.float32).reshape(n, 3).contiguous()
        rt.restart_F(ctx, model, x.contiguous(), target, mpm_state.particle_F_trial)
        x.copy_(target)
