"""One solver configuration ("mpm_config") for the MPM path.

The reference spreads a simulation's settings over Python literals: the
WarpMPMWrapper constructor (grid 100, grid_lim 1, frame_dt 0.01, sim_dt
1e-3, gravity (0, 0, -9.8); warp_mpm_wrapper.py:92-110), the keys that
MPMWARPDiff.set_parameters_dict reads (mpm_solver_diff.py:38-104) and the
collider registration calls (mpm_solver_diff.py:437-938).  Here they are one
dict, or a YAML or JSON file read with safe loaders:

    grid_size: 64
    grid_lim: 1.0
    frame_dt: 0.01
    sim_dt: 5.0e-5
    gravity: [0.0, 0.0, -9.8]
    material: neo-hookean-von-mises     # a name of MATERIALS or its id
    yield_stress: 4000.0
    grid_v_damping_scale: 0.98
    colliders:
      - {type: surface, point: [0, 0, 0.1], normal: [0, 0, 1], surface: sticky}
      - {type: geometry, shape: sphere, center: [0.5, 0.5, 0.5], radius: 0.1}

load_mpm_config validates and fills defaults; solver_parameters gives the
set_parameters_dict subset; apply_to_solver and make_wrapper hand a
configuration to the drop-in API.  Unknown keys are errors (a misspelt key
would otherwise be ignored silently, as the reference's dict lookups do).
"""
from __future__ import annotations

import copy
import json
import os

import yaml

from empm_amd.diff_simulator.warp_mpm.mpm_solver_diff import MATERIALS

# set_parameters_dict's keys (mpm_solver_diff.py:38-104)
SOLVER_KEYS = ("material", "yield_stress", "hardening", "xi", "friction_angle", "rpic_damping", "plastic_viscosity",
               "softening", "grid_v_damping_scale")

DEFAULTS = {
    "grid_size": 100,
    "grid_lim": 1.0,
    "frame_dt": 0.01,
    "sim_dt": 1e-3,
    "gravity": [0.0, 0.0, -9.8],
    "colliders": [],
}

# collider type -> (the solver method it registers through, required keys)
COLLIDERS = {
    "surface": ("add_surface_collider", ("point", "normal")),
    "cuboid": ("set_velocity_on_cuboid", ("point", "size", "velocity")),
    "bounding_box": ("add_bounding_box", ()),
    "geometry": ("add_geometry_collider", ("shape",)),
    "geometry_mover": ("set_grid_velocity_on_geometry", ("shape",)),
}

_NUMBER = (int, float)


def _vec3(name, v):
    if not (isinstance(v, (list, tuple)) and len(v) == 3 and all(isinstance(a, _NUMBER) for a in v)):
        raise ValueError(f"mpm_config: {name} must be three numbers, got {v!r}")
    return [float(a) for a in v]


def _read(src):
    if isinstance(src, dict):
        return copy.deepcopy(src)
    path = os.fspath(src)
    with open(path) as f:
        if path.endswith(".json"):
            return json.load(f)
        return yaml.safe_load(f) or {}


def load_mpm_config(src) -> dict:
    """A validated configuration with defaults filled in, from a dict or a
    YAML / JSON file path.  Raises ValueError on unknown keys, bad values or
    an unknown material, naming the key."""
    raw = _read(src)
    if not isinstance(raw, dict):
        raise ValueError("mpm_config: the configuration must be a mapping")
    unknown = sorted(set(raw) - set(DEFAULTS) - set(SOLVER_KEYS))
    if unknown:
        raise ValueError(f"mpm_config: unknown keys {unknown}")
    cfg = copy.deepcopy(DEFAULTS)
    cfg.update(raw)
    if not isinstance(cfg["grid_size"], int) or not 4 <= cfg["grid_size"] <= 1024:
        raise ValueError(f"mpm_config: grid_size must be an integer in [4, 1024], got {cfg['grid_size']!r}")
    for k in ("grid_lim", "frame_dt", "sim_dt"):
        if not isinstance(cfg[k], _NUMBER) or not cfg[k] > 0:
            raise ValueError(f"mpm_config: {k} must be a positive number, got {cfg[k]!r}")
        cfg[k] = float(cfg[k])
    cfg["gravity"] = _vec3("gravity", cfg["gravity"])
    if "material" in cfg:
        m = cfg["material"]
        if isinstance(m, int) and not isinstance(m, bool):
            names = [k for k, v in MATERIALS.items() if v == m]
            if not names:
                raise ValueError(f"mpm_config: material id {m} outside {sorted(MATERIALS.values())}")
            cfg["material"] = names[0]
        elif m not in MATERIALS:
            raise ValueError(f"mpm_config: unknown material {m!r} (one of {sorted(MATERIALS)})")
    for k in SOLVER_KEYS:
        if k != "material" and k in cfg:
            if not isinstance(cfg[k], _NUMBER) or isinstance(cfg[k], bool):
                raise ValueError(f"mpm_config: {k} must be a number, got {cfg[k]!r}")
            cfg[k] = float(cfg[k])
    if not isinstance(cfg["colliders"], list):
        raise ValueError("mpm_config: colliders must be a list")
    cols = []
    for i, c in enumerate(cfg["colliders"]):
        if not isinstance(c, dict) or c.get("type") not in COLLIDERS:
            raise ValueError(f"mpm_config: collider {i} needs a type in {sorted(COLLIDERS)}")
        missing = [k for k in COLLIDERS[c["type"]][1] if k not in c]
        if missing:
            raise ValueError(f"mpm_config: collider {i} ({c['type']}) misses {missing}")
        cols.append(copy.deepcopy(c))
    cfg["colliders"] = cols
    return cfg


def solver_parameters(cfg: dict) -> dict:
    """The keys of cfg that MPMWARPDiff.set_parameters_dict reads."""
    return {k: cfg[k] for k in SOLVER_KEYS if k in cfg}


def apply_to_solver(solver, mpm_model, mpm_state, cfg: dict, device="cuda:0"):
    """On an initialized solver (MPMWARPDiff.initialize, which starts its
    collider list): set_parameters_dict with cfg's solver keys, the model's
    gravity (gravitational_accelaration, the reference's spelling), then
    every collider of cfg registered in order (the order fixes each
    collider's index, as in the reference)."""
    solver.set_parameters_dict(mpm_model, mpm_state, solver_parameters(cfg), device=device)
    mpm_model.gravitational_accelaration = tuple(cfg["gravity"])
    for c in cfg["colliders"]:
        method, _ = COLLIDERS[c["type"]]
        args = {k: v for k, v in c.items() if k != "type"}
        if c["type"] in ("geometry", "geometry_mover"):
            getattr(solver, method)(args)
        else:
            getattr(solver, method)(**args)


def make_wrapper(cfg: dict, device="cuda:0", requires_grad=False, **callbacks):
    """A WarpMPMWrapper built from cfg's grid, time steps and gravity; its
    material parameters (material_params) take cfg's solver keys, which
    initialize() passes to set_parameters_dict.  Colliders stay to the
    caller's add_* calls, whose arguments differ between the wrapper and the
    solver (warp_mpm_wrapper.py:194-209)."""
    from empm_amd.diff_simulator.warp_mpm.warp_mpm_wrapper import WarpMPMWrapper

    w = WarpMPMWrapper(grid_size=cfg["grid_size"], grid_lim=cfg["grid_lim"], frame_dt=cfg["frame_dt"],
                       sim_dt=cfg["sim_dt"], device=device, requires_grad=requires_grad, **callbacks)
    w.gravity = list(cfg["gravity"])
    params = dict(w.material_params)
    params.update(solver_parameters(cfg))
    w.material_params = params
    return w
